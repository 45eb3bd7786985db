#!/usr/bin/env python3
"""Throughput of the MiniFS solver path on one MI355X.

Default workload: the box + Gaussian-bump mesh at 100x100x100 hexes x 6
tetrahedra (6.0M cells, 12.06M faces), Mach 0.5, AoA 0, CFL 10, plain FP64,
TDTree-ordered.  One step is one iteration of the reference driver loop
(src/main.cpp:201-291): residual (Roe flux), local time step, LU-SGS
forward/backward sweeps with the flow-field update, Green-Gauss gradients.
The mesh and state are resident in HBM before the timed region.

Prints one JSON line.  `roofline` describes the kernel with the largest
share of the timed region, timed with HIP events on the solver stream (the
timed region brackets the two sweeps only; `kernel_ms_per_step` times every
kernel group over further iterations);
`cpu_baseline` times the reference's own OpenMP build (oracle/_ref) on all
host threads for a few iterations of the same mesh and flow, with its
serial build beside it; `mixed` repeats the
timed region with the mixed-precision context (FP32 implicit operator,
FP64 residual and update) on the same mesh.

Other modes (one JSON line each):
  --mode converge  both precisions run to a residual drop on the same mesh
                   and flow (default Mach 0.8, AoA 1.25); errors of the
                   mixed run against the FP64 run
  --mode batch     many independent solves of random size and flow stepped
                   together (mfs_batch_*), aggregate throughput
  --mode serve     Poisson arrivals of solves of random size, served by the
                   batching server against one-at-a-time plain FP64 solves
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
    # backward sweep + update: as the forward sweep plus the state update
    "bwd_sweep": (40, 264, 40, 0),
    # gradient of one iteration + residual/time-step pass of the next (mfs_step,
    # tetrahedra): the union of the two: face tables (flux_ts's 56 + the 3
    # face nodes), q + vol read once, res, dt, Diag and 15 gradient outputs
    # written, ghosts, node values
    "grad_flux_ts": (68, 184, 80, 40),
    # node values: 4 weighted entries per tetrahedron (index + weight), q, node write
    "node": (0, 88, 76, 52),
    # Green-Gauss: face nodes + geometry, q, node values, vol, 15 gradient outputs
    "gradient": (52, 128, 80, 40),
}
# kernels that mfs_step runs on the context's second stream by default
SIDE_KERNELS = ("node", "gradient")
# a sweep run as two launches (wide levels, thin tail) is timed as a whole
# and per launch; each launch's algorithmic bytes are its cells' share of
# the sweep's
SWEEP_LAUNCHES = {"fwd_sweep": ("fwd_sweep_wide", "fwd_sweep_tail"), "bwd_sweep": ("bwd_sweep_tail", "bwd_sweep_wide")}
# iterations after the timed region with every kernel group timed
KERNEL_TABLE_STEPS = 10
# whole iteration, plain FP64 (268 nF + 840 nC + 356 nB + 92 nN)
ITER_BYTES = (268, 840, 356, 92)


# A shallower TDTree cell order than the reference's 4 parts / 512-cell
# leaves: on the 100^3 x 6 mesh 82 sweep levels instead of 342.
SHALLOW_TREE = (2, 128)


def alg_bytes(coef, nF, nC, nB, nN):
    return coef[0] * nF + coef[1] * nC + coef[2] * nB + coef[3] * nN


def host_cpu() -> dict:
    """The host's CPU model, logical CPU count and the OpenMP binding in use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND")}


def _run_reference(exe: str, mesh: str, bc: str, iters: int, mach: float, alpha: float, cfl: float,
                   env: dict | None = None) -> tuple[float, float]:
    """Runs a reference build (oracle/_ref) for `iters` iterations on a saved
    mesh; returns (loop time, whole-run wall time) in seconds."""
    with tempfile.TemporaryDirectory() as out:
        t0 = time.perf_counter()
        subprocess.run([exe, mesh, bc, out, repr(mach), repr(alpha), repr(cfl), str(iters), "", "nogeo"], check=True,
                       stdout=subprocess.DEVNULL, timeout=1200, env=env)
        wall = time.perf_counter() - t0
        with open(os.path.join(out, "summary.txt")) as f:
            first = f.readline().split()
    kv = dict(zip(first[0::2], first[1::2]))
    return float(kv["t_loop"]), wall


def cpu_samples(mesh: str, bc: str, nC: int, label: str, args) -> tuple[dict | None, dict | None]:
    """The CPU path beside the GPU, on the same saved mesh.

    cpu_baseline: the reference's own OpenMP face-colouring build (command:22:
    -fopenmp -DMF_OPENMP -DOMP_CellColor -DOMP_FaceColor, built from the
    reference sources into oracle/_ref/ref_driver_omp) on all host threads
    with OMP_PROC_BIND=close, `cpu_iters` iterations, loop time only
    (src/main.cpp:183-196, 353-354).  Its LU-SGS runs over cell colours, so
    its sums differ from the serial order: a timing baseline, not a parity
    reference.
    The serial reference build (oracle/_ref/ref_driver) is timed beside it on
    the same mesh for `cpu_serial_iters` iterations."""
    omp = os.path.join(ROOT, "oracle", "_ref", "ref_driver_omp")
    ser = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    base = serial = None
    if os.path.exists(omp) and args.cpu_iters > 0:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        t_loop, wall = _run_reference(omp, mesh, bc, args.cpu_iters, args.mach, args.alpha, args.cfl, env)
        host = host_cpu()
        host["omp_proc_bind"] = "close"
        base = {
            "value": round(nC * args.cpu_iters / t_loop / 1e6, 4),
            "unit": "M cell-iters/s",
            "cores": threads,
            "kind": "reference",
            "sample": f"reference OpenMP face-colouring build (oracle/_ref/ref_driver_omp: g++ -O3 -fopenmp "
                      f"-DMF_OPENMP -DOMP_CellColor -DOMP_FaceColor), {threads} threads, OMP_PROC_BIND=close, "
                      f"{args.cpu_iters} iterations of the {label} = {nC} cell mesh, same flow condition, loop "
                      f"time only ({t_loop:.1f} s of {wall:.1f} s incl. mesh read and set-up)",
            "host": host,
        }
    if os.path.exists(ser) and args.cpu_serial_iters > 0:
        t_loop, wall = _run_reference(ser, mesh, bc, args.cpu_serial_iters, args.mach, args.alpha, args.cfl)
        serial = {
            "value": round(nC * args.cpu_serial_iters / t_loop / 1e6, 4),
            "unit": "M cell-iters/s",
            "cores": 1,
            "kind": "reference",
            "sample": f"reference serial build (oracle/_ref/ref_driver, g++ -O3 -ffp-contract=off), "
                      f"{args.cpu_serial_iters} iterations of the same mesh, loop time only ({t_loop:.1f} s)",
        }
    return base, serial


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


def timed_steps(s, steps: int, warmup: int, kernel_timing: int):
    """Warm-up steps, then `steps` timed steps bracketed by synchronisation;
    returns wall seconds and the per-kernel event times (kernel_timing: 0
    none, 1 every kernel group, 2 the two sweeps only)."""
    s.step(warmup)
    s.sync()
    s.enable_timing(kernel_timing)
    s.reset_timing()
    t1 = time.perf_counter()
    s.step(steps)
    s.sync()
    elapsed = time.perf_counter() - t1
    kern = {}
    for k in list(KERNEL_BYTES) + [part for parts in SWEEP_LAUNCHES.values() for part in parts]:
        ms, cnt = s.timing(k)
        if cnt:
            kern[k] = (ms, cnt)
    s.enable_timing(False)
    return elapsed, kern


def mode_fp64(args, cpu_dir: str | None = None) -> dict:
    import minifs_amd as mfs

    t0 = time.time()
    p = mfs.make_params(mach=args.mach, alpha=args.alpha, cfl=args.cfl)
    g = mfs.Grid.box_bump(args.n, seed=args.seed).decouple(*args.tree)
    if cpu_dir:  # the CPU path reads the same mesh (the reference's FlowStar format)
        g.save(os.path.join(cpu_dir, "m.mfl"), os.path.join(cpu_dir, "m.bc"))
    g.preprocess(p)
    info = g.info
    s = mfs.Solver(g, p)
    setup_s = time.time() - t0
    nC, nF, nB, nN = info["nTCell"], info["nTFace"], info["nBFace"], info["nTNode"]
    # the timed region brackets only the sweeps with HIP events (each event
    # record idles the stream a few us); every kernel's time comes from
    # further iterations with all kernels timed
    elapsed, kern = timed_steps(s, args.steps, args.warmup, 0 if args.no_kernel_timing else 2)
    kern_all = {} if args.no_kernel_timing else timed_steps(s, KERNEL_TABLE_STEPS, 0, 1)[1]
    ms_per_step = elapsed / args.steps * 1e3
    value = nC * args.steps / elapsed / 1e6
    hist = s.history()
    info_ctx = s.info()
    tail_cells = info_ctx["sweep_tail_cells"]

    def launch_bytes(k):
        """Algorithmic bytes of one launch of kernel group k."""
        for sweep, parts in SWEEP_LAUNCHES.items():
            if k in parts:
                cells = tail_cells if k.endswith("_tail") else nC - tail_cells
                return alg_bytes(KERNEL_BYTES[sweep], nF, nC, nB, nN) * cells / nC
        return alg_bytes(KERNEL_BYTES[k], nF, nC, nB, nN)

    roof = None
    sweeps = None
    if kern:
        # the dominant kernel launch of the critical path: the solver
        # stream's launches (node values and the gradient run on the
        # context's second stream beside them); a split sweep counts as its
        # two launches
        split = [k for k in SWEEP_LAUNCHES if SWEEP_LAUNCHES[k][0] in kern]
        launches = [k for k in kern if k not in SIDE_KERNELS and k not in split]
        dom = max(launches, key=lambda k: kern[k][0])
        ms, cnt = kern[dom]
        avg_s = ms / cnt / 1e3
        b = launch_bytes(dom)
        achieved = b / avg_s / 1e9
        sweeps = {}
        for k in SWEEP_LAUNCHES:
            if k not in kern:
                continue
            sb = alg_bytes(KERNEL_BYTES[k], nF, nC, nB, nN)
            span = kern[k][0] / kern[k][1]
            sweeps[k] = {"span_ms": round(span, 4), "alg_bytes": int(sb),
                         "achieved_GBs": round(sb / (span / 1e3) / 1e9, 2)}
            for part in SWEEP_LAUNCHES[k]:
                if part in kern:
                    pm = kern[part][0] / kern[part][1]
                    sweeps[k][part + "_ms"] = round(pm, 4)
                    sweeps[k][part + "_achieved_GBs"] = round(launch_bytes(part) / (pm / 1e3) / 1e9, 2)
        traffic = None
        tf = os.path.join(ROOT, "profiles", f"traffic_{dom}.json")
        if os.path.exists(tf):
            with open(tf) as f:
                traffic = json.load(f).get("bytes_per_launch")
        # the sweeps move HBM bytes, but their time is set by the DAG depth
        # (sweep levels x a dependency hop and a chunk's arithmetic), not by
        # bandwidth
        limiter = ("dependency waits and scattered gathers (wide sweep levels)" if dom.endswith("_wide") else
                   "dependency latency (thin sweep levels)" if dom.endswith("_tail") else
                   "dependency latency (sweep levels)" if dom.endswith("sweep") else "hbm")
        roof = {"kernel": dom, "bound": "hbm", "limiter": limiter, "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "alg_bytes_per_launch": int(b), "avg_launch_ms": round(ms / cnt, 4)}
        if dom.endswith("_wide"):
            roof["launch"] = (f"the sweep's wide levels: {nC - tail_cells} of its {nC} cells; its last "
                              f"{tail_cells} cells (the thin tail) run as a second launch")
    iter_b = alg_bytes(ITER_BYTES, nF, nC, nB, nN)
    s.close()
    # the same iteration with the gradient on the solver stream (no second
    # stream beside the sweeps): each kernel's own duration, and what the
    # overlap gains
    isolated = None
    if kern and not args.no_kernel_timing:
        s1 = mfs.Solver(g, p, gradient_stream=1)
        el1, kern1 = timed_steps(s1, args.steps, args.warmup, 1)
        s1.close()
        dom = roof["kernel"]
        ms1, cnt1 = kern1[dom]
        b = launch_bytes(dom)
        a1 = b / (ms1 / cnt1 / 1e3) / 1e9
        isolated = {"gradient_stream": 1, "ms_per_step": round(el1 / args.steps * 1e3, 4),
                    "value": round(nC * args.steps / el1 / 1e6, 3),
                    "roofline": {"kernel": dom, "achieved": round(a1, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(a1 / HBM_PEAK_GBS, 5), "avg_launch_ms": round(ms1 / cnt1, 4)},
                    "kernel_ms_per_step": {k: round(v[0] / v[1], 4) for k, v in kern1.items()}}
    mixed = None
    if not args.no_mixed:
        sm = mfs.Solver(g, p, precision=mfs.MFS_MIXED)
        el_m, _ = timed_steps(sm, args.steps, args.warmup, 0 if args.no_kernel_timing else 2)
        kern_m = {} if args.no_kernel_timing else timed_steps(sm, KERNEL_TABLE_STEPS, 0, 1)[1]
        mixed = {"value": round(nC * args.steps / el_m / 1e6, 3), "unit": "M cell-iters/s",
                 "ms_per_step": round(el_m / args.steps * 1e3, 4), "speedup_vs_fp64": round(elapsed / el_m, 3),
                 "dtype": "f32 operator / f64 residual+update",
                 "kernel_ms_per_step": {k: round(v[0] / v[1], 4) for k, v in kern_m.items()}}
        sm.close()
    g.close()
    shallow = None
    if not args.no_shallow and tuple(args.tree) != SHALLOW_TREE:
        # the same mesh and flow under a shallower cell order (a set-up
        # choice: any order is a valid input to the path)
        g2 = mfs.Grid.box_bump(args.n, seed=args.seed).decouple(*SHALLOW_TREE).preprocess(p)
        shallow = {"ordering": f"TDTree {SHALLOW_TREE[0]} parts, {SHALLOW_TREE[1]}-cell leaves"}
        for prec, key in ((mfs.MFS_FP64, "fp64"), (mfs.MFS_MIXED, "mixed")):
            s2 = mfs.Solver(g2, p, precision=prec)
            if prec == mfs.MFS_FP64:
                shallow["sweep_levels"] = s2.info()["sweep_levels"]
            el2, _ = timed_steps(s2, args.steps, args.warmup, 0)
            shallow[key] = {"value": round(nC * args.steps / el2 / 1e6, 3), "unit": "M cell-iters/s",
                            "ms_per_step": round(el2 / args.steps * 1e3, 4)}
            s2.close()
        g2.close()
    return {
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
                   "cfl": args.cfl, "seed": args.seed, "ordering": f"TDTree {args.tree[0]} parts, {args.tree[1]}-cell leaves",
                   "sweep_levels": info_ctx["sweep_levels"]},
        "roofline": roof,
        "iteration_roofline": {"alg_bytes_per_iter": int(iter_b),
                               "achieved_GBs": round(iter_b / (ms_per_step / 1e3) / 1e9, 2),
                               "frac": round(iter_b / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 5)},
        "sweeps": sweeps,
        "kernel_ms_per_step": {k: round(v[0] / v[1], 4) for k, v in kern_all.items()},
        "kernel_table": f"HIP events around every kernel group, {KERNEL_TABLE_STEPS} iterations after the timed "
                        "region (whose events bracket the sweeps only)",
        "gradient_on_solver_stream": isolated,
        "mixed": mixed,
        "shallow_order": shallow,
        "setup_s": round(setup_s, 2),
        "last_residual": [float(x) for x in hist[-1][:5]] if len(hist) else None,
        "cpu_baseline": None,
        "cpu_reference_serial": None,
    }


def mode_converge(args) -> dict:
    """Both precisions to a residual drop on one mesh: iterations, time, and
    the mixed run's state and wall-force errors against the FP64 run."""
    import numpy as np

    import minifs_amd as mfs

    p = mfs.make_params(mach=args.mach, alpha=args.alpha, cfl=args.cfl)
    g = mfs.Grid.box_bump(args.n, seed=args.seed).decouple(*args.tree).preprocess(p)
    nC = g.info["nTCell"]
    runs = {}
    for prec, name in ((mfs.MFS_FP64, "fp64"), (mfs.MFS_MIXED, "mixed")):
        s = mfs.Solver(g, p, precision=prec)
        t1 = time.perf_counter()
        done = 0
        while done < args.max_iters:
            k = min(args.check_every, args.max_iters - done)
            s.step(k)
            done += k
            h = s.history()
            if done % (20 * args.check_every) == 0:
                print(f"[converge] {name}: {done} iterations, drop {h[-1, 0] / h[0, 0]:.3e}, "
                      f"{time.perf_counter() - t1:.1f} s", file=sys.stderr, flush=True)
            if h[-1, 0] <= args.drop * h[0, 0]:
                break
        s.sync()
        el = time.perf_counter() - t1
        st = {nm: s.field(nm) for nm in ("rho", "u", "v", "w", "p")}
        F, lift = mfs.wall_force(g, st["p"], p)
        runs[name] = {"iters": len(h), "seconds": round(el, 3), "drop": float(h[-1, 0] / h[0, 0]),
                      "M_cell_iters_per_s": round(nC * len(h) / el / 1e6, 3), "lift": lift, "state": st}
        print(f"[converge] {name}: {len(h)} iterations, {el:.1f} s, drop {h[-1, 0] / h[0, 0]:.3e}", file=sys.stderr,
              flush=True)
        s.close()
    err = {}
    for nm in ("rho", "u", "v", "w", "p"):
        a, b = runs["fp64"]["state"][nm], runs["mixed"]["state"][nm]
        scale = (p.p_bar if nm == "p" else float(np.max(np.abs(a))))
        err[nm] = {"l2_rel": float(np.linalg.norm(a - b) / max(np.linalg.norm(a), 1e-300)),
                   "linf_rel": float(np.max(np.abs(a - b)) / scale)}
    for r in runs.values():
        r.pop("state")
    g.close()
    return {"metric": "converged mixed vs FP64 solve", "mode": "converge",
            "config": {"workload": f"box+bump {args.n}^3x6 tets", "cells": nC, "mach": args.mach,
                       "alpha_deg": args.alpha, "cfl": args.cfl, "drop": args.drop},
            "fp64": runs["fp64"], "mixed": runs["mixed"],
            "speedup_time_to_converge": round(runs["fp64"]["seconds"] / runs["mixed"]["seconds"], 3),
            "state_error": err,
            "lift_rel_error": abs(runs["mixed"]["lift"] - runs["fp64"]["lift"]) / abs(runs["fp64"]["lift"])}


def mode_batch_converge(args) -> dict:
    """The same seeded solves, all present at the start, each run to a
    residual drop of --batch-drop in one device batch (converged solves
    leave it), against the same solves run one at a time on the plain FP64
    path."""
    import minifs_amd as mfs
    from minifs_amd.serve import Server, make_batch_requests, run_sequential

    def reqs():
        return make_batch_requests(args.solves, args.batch_lo, args.batch_hi, args.batch_seed)

    prec = mfs.MFS_MIXED if args.precision == "mixed" else mfs.MFS_FP64
    batch_reqs = reqs()
    srv = Server(prec, hbm_budget=args.hbm_budget, drop=args.batch_drop, max_iters=args.max_iters,
                 check_every=args.check_every, workers=min(16, os.cpu_count() or 4), tree=args.tree)
    served = srv.run(batch_reqs, progress=lambda r: print(f"[batch] solve {r.rid} {r.n_cells} cells {r.iters} it "
                                                          f"{r.t_done:.2f} s", file=sys.stderr, flush=True))
    print(json.dumps({"batched": served}), file=sys.stderr, flush=True)
    out = {"metric": "batched solves to a residual drop", "mode": "batch", "batched": served,
           "config": {"solves": args.solves, "hex_per_axis": [args.batch_lo, args.batch_hi], "seed": args.batch_seed,
                      "precision": args.precision, "drop": args.batch_drop, "check_every": args.check_every,
                      "tree": list(args.tree)}}
    if not args.no_sequential:
        seq_reqs = reqs()
        seq = run_sequential(seq_reqs, drop=args.batch_drop, max_iters=args.max_iters, check_every=args.check_every,
                             tree=args.tree,
                             progress=lambda r: print(f"[sequential] solve {r.rid} {r.n_cells} cells {r.iters} it "
                                                      f"{r.t_done:.2f} s", file=sys.stderr, flush=True))
        out["sequential_fp64"] = seq
        out["wall_speedup"] = round(seq["wall_s"] / served["wall_s"], 3)
    return out


def mode_batch(args) -> dict:
    """Independent solves with nx,ny,nz ~ U[lo,hi], bump ~ U[.02,.08], Mach ~
    U[.3,.85], AoA ~ U[-3,5] (seeded), stepped together."""
    import concurrent.futures as cf

    import numpy as np

    import minifs_amd as mfs

    rng = np.random.default_rng(args.batch_seed)
    specs = [(int(rng.integers(args.batch_lo, args.batch_hi + 1)), int(rng.integers(args.batch_lo, args.batch_hi + 1)),
              int(rng.integers(args.batch_lo, args.batch_hi + 1)), float(rng.uniform(0.02, 0.08)),
              float(rng.uniform(0.3, 0.85)), float(rng.uniform(-3.0, 5.0)), int(rng.integers(1, 2**31)))
             for _ in range(args.solves)]
    prec = mfs.MFS_MIXED if args.precision == "mixed" else mfs.MFS_FP64

    def build(sp):
        nx, ny, nz, bh, mach, alpha, seed = sp
        p = mfs.make_params(mach=mach, alpha=alpha, cfl=args.cfl)
        return mfs.Grid.box_bump(nx, ny, nz, bump_height=bh, seed=seed).decouple(*args.tree).preprocess(p), p

    t0 = time.time()
    with cf.ThreadPoolExecutor(min(16, os.cpu_count() or 4)) as pool:
        built = list(pool.map(build, specs))
    setup_s = time.time() - t0
    from minifs_amd.serve import tuned_solver

    b = mfs.Batch(prec)
    members = [b.attach(tuned_solver(g, p, prec)) for g, p in built]
    cells = sum(g.info["nTCell"] for g, _ in built)
    b.step(args.warmup)
    b.sync()
    t1 = time.perf_counter()
    b.step(args.steps)
    b.sync()
    el = time.perf_counter() - t1
    out = {"metric": "aggregate M cell-iters/s of a batch", "mode": "batch", "value": round(cells * args.steps / el / 1e6, 3),
           "unit": "M cell-iters/s", "ms_per_step": round(el / args.steps * 1e3, 4), "steps": args.steps,
           "config": {"solves": args.solves, "cells_total": cells, "hex_per_axis": [args.batch_lo, args.batch_hi],
                      "seed": args.batch_seed, "precision": args.precision}, "setup_s": round(setup_s, 2)}
    # the same solves one at a time (same precision) for the batching gain
    if not args.no_sequential:
        t2 = time.perf_counter()
        for g, p in built:
            s = tuned_solver(g, p, prec)
            s.step(args.steps)
            s.sync()
            s.close()
        out["sequential_value"] = round(cells * args.steps / (time.perf_counter() - t2) / 1e6, 3)
    b.close()
    for g, _ in built:
        g.close()
    return out


def mode_serve(args) -> dict:
    import minifs_amd as mfs
    from minifs_amd.serve import Server, make_requests, run_sequential

    prec = mfs.MFS_MIXED if args.precision == "mixed" else mfs.MFS_FP64
    reqs = make_requests(args.requests, args.rate, args.cells_lo, args.cells_hi, args.serve_seed)
    srv = Server(prec, hbm_budget=args.hbm_budget, drop=args.serve_drop, max_iters=args.max_iters,
                 check_every=args.check_every, workers=min(16, os.cpu_count() or 4), tree=args.tree)
    served = srv.run(reqs, progress=lambda r: print(f"[serve] request {r.rid} {r.n_cells} cells {r.iters} it "
                                                    f"{r.latency:.2f} s", file=sys.stderr, flush=True))
    print(json.dumps({"served": served}), file=sys.stderr, flush=True)
    base = None
    if not args.no_sequential:
        again = make_requests(args.requests, args.rate, args.cells_lo, args.cells_hi, args.serve_seed)
        base = run_sequential(again, drop=args.serve_drop, max_iters=args.max_iters, check_every=args.check_every,
                              tree=args.tree,
                              progress=lambda r: print(f"[sequential] request {r.rid} {r.n_cells} cells {r.iters} it "
                                                       f"{r.latency:.2f} s", file=sys.stderr, flush=True))
        same = all(a.state_sha256 == b.state_sha256 and a.iters == b.iters
                   for a, b in zip(sorted(reqs, key=lambda r: r.rid), sorted(again, key=lambda r: r.rid))) \
            if prec == mfs.MFS_FP64 else None
        base["results_equal_to_served"] = same
    return {"metric": "served solves/hr + p50/p99 time-to-converge", "mode": "serve", "served": served,
            "sequential_fp64": base,
            "solves_per_hour_gain": round(served["solves_per_hour"] / base["solves_per_hour"], 3) if base else None,
            "config": {"requests": args.requests, "rate_per_s": args.rate, "cells": [args.cells_lo, args.cells_hi],
                       "drop": args.serve_drop, "precision": args.precision, "hbm_budget": args.hbm_budget,
                       "seed": args.serve_seed, "tree": list(args.tree)}}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--mode", choices=["fp64", "converge", "batch", "serve"], default="fp64")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=100, help="hexes per axis (cells = 6 n^3)")
    ap.add_argument("--mach", type=float, default=None, help="default 0.5 (0.8 for --mode converge)")
    ap.add_argument("--alpha", type=float, default=None, help="default 0 (1.25 for --mode converge)")
    ap.add_argument("--cfl", type=float, default=10.0)
    ap.add_argument("--seed", type=int, default=20260101)
    ap.add_argument("--tree", type=lambda t: tuple(int(x) for x in t.split(",")), default=(4, 512),
                    metavar="PARTS,LEAF", help="TDTree parameters of the mesh order (the reference's: 4,512)")
    ap.add_argument("--cpu-iters", type=int, default=10,
                    help="iterations of the reference's OpenMP build on the bench mesh (0 = skip)")
    ap.add_argument("--cpu-serial-iters", type=int, default=2,
                    help="iterations of the reference's serial build on the bench mesh (0 = skip)")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision repeat of the timed region")
    ap.add_argument("--no-shallow", action="store_true", help="skip the repeat on the shallower cell order")
    ap.add_argument("--drop", type=float, default=1e-10, help="converge mode: residual drop")
    ap.add_argument("--max-iters", type=int, default=20000)
    ap.add_argument("--check-every", type=int, default=50)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed", help="batch and serve modes")
    ap.add_argument("--solves", type=int, default=64)
    ap.add_argument("--batch-lo", type=int, default=15)
    ap.add_argument("--batch-hi", type=int, default=60)
    ap.add_argument("--batch-seed", type=int, default=4001)
    ap.add_argument("--batch-drop", type=float, default=1e-8,
                    help="batch mode: run each solve to this residual drop (0: a fixed number of steps)")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--cells-lo", type=float, default=20000)
    ap.add_argument("--cells-hi", type=float, default=3000000)
    ap.add_argument("--serve-drop", type=float, default=1e-8)
    ap.add_argument("--serve-seed", type=int, default=5001)
    ap.add_argument("--hbm-budget", type=float, default=200e9)
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (for scaling rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration, 5)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2
    conv = args.mode == "converge"
    if args.mach is None:
        args.mach = 0.8 if conv else 0.5
    if args.alpha is None:
        args.alpha = 1.25 if conv else 0.0
    rank = int(os.environ.get("RANK", "0"))

    if args.mode == "converge":
        line = mode_converge(args)
    elif args.mode == "batch":
        line = mode_batch_converge(args) if args.batch_drop > 0 else mode_batch(args)
    elif args.mode == "serve":
        line = mode_serve(args)
    else:
        cpu = rank == 0 and (args.cpu_iters > 0 or args.cpu_serial_iters > 0)
        with tempfile.TemporaryDirectory() as td:
            line = mode_fp64(args, td if cpu else None)
            if cpu:
                try:
                    line["cpu_baseline"], line["cpu_reference_serial"] = cpu_samples(
                        os.path.join(td, "m.mfl"), os.path.join(td, "m.bc"), line["config"]["cells"],
                        f"{args.n}^3x6 box+bump", args)
                except Exception as e:  # reported, not fatal: the GPU number stands on its own
                    line["cpu_baseline"] = {"error": str(e)}
    if rank == 0:
        print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
